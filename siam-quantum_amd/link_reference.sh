#!/bin/bash
# Links the reference program against libsqamd.so as a drop-in.
#
#   siam-quantum_amd/link_reference.sh <reference src dir> <work dir> [extra link flags]
#
# Every reference source file is compiled as it is; in each object, the
# global functions that libsqamd.so also defines (the J/K engine, uhf/rhf,
# the density DIIS, the eigen-solvers, the int.h/fgamma.h scalar functions,
# the readers and checkpoint functions) are renamed to sqref_<name> with
# objcopy, so that every call to them from the remaining reference code
# binds to the library, and the library's copy is the one that runs.  The
# program is then linked against the library.  Extra link flags (the
# reference's LAPACK, e.g. "-llapack -lrefblas -lgfortran") go last; without
# them the link reports the symbols left unresolved and stops with status 3
# if any is outside LAPACK (dspgvd_, dspgvx_, dgesv_), which only the
# renamed, unused eigen-solver copies call.
set -euo pipefail
SRC=${1:?reference src dir}
WORK=${2:?work dir}
shift 2
HERE=$(cd "$(dirname "$0")" && pwd)
LIB=$HERE/lib/libsqamd.so
[ -f "$LIB" ] || { echo "link_reference: $LIB is not built" >&2; exit 2; }
mkdir -p "$WORK"
cd "$WORK"
for f in "$SRC"/*.c; do
  gcc -O2 -w -c "$f" -o "$(basename "$f" .c).o"
done
nm -D --defined-only "$LIB" | awk '$2=="T"{print $3}' | sort -u > libsqamd.syms
for o in *.o; do
  args=()
  for s in $(nm --defined-only -g "$o" | awk '{print $3}' | sort -u | comm -12 - libsqamd.syms); do
    args+=(--redefine-sym "$s=sqref_$s")
  done
  if [ ${#args[@]} -gt 0 ]; then objcopy "${args[@]}" "$o"; fi
done
if [ $# -gt 0 ]; then
  gcc -o sq ./*.o -L"$HERE/lib" -lsqamd -Wl,-rpath,"$HERE/lib" -lm "$@"
  echo "link_reference: linked $WORK/sq"
  exit 0
fi
gcc -o sq ./*.o -L"$HERE/lib" -lsqamd -Wl,-rpath,"$HERE/lib" -lm -Wl,--warn-unresolved-symbols 2> link.log || true
grep -o "undefined reference to \`[^']*'" link.log | sed "s/undefined reference to \`//; s/'//" | sort -u > unresolved.txt || true
echo "link_reference: unresolved: $(tr '\n' ' ' < unresolved.txt)"
if grep -v -x -e dspgvd_ -e dspgvx_ -e dgesv_ unresolved.txt; then
  echo "link_reference: symbols other than LAPACK are unresolved" >&2
  exit 3
fi
echo "link_reference: linked $WORK/sq (LAPACK left to the reference's own build)"
