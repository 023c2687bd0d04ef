#!/bin/sh
# Builds oracle/_ref/libref.so from the reference's CPU sources where they lie
# (REF_SRC, default /root/reference/src) plus oracle/ref_harness.cpp, with the
# reference makefile's CPU flags (makefile:3: -std=c++11 -fopenmp -O2).
# Test infrastructure only; the output stays out of git (.gitignore: /oracle/_ref/).
# Exits 0 without building when the reference sources are absent.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
REF_SRC=${REF_SRC:-/root/reference/src}
if [ ! -f "$REF_SRC/cpu_3d.cpp" ]; then
  echo "build_ref: reference sources not found at $REF_SRC; skipping oracle/_ref"
  exit 0
fi
OUT="$HERE/_ref"
mkdir -p "$OUT"
CXX=${CXX:-g++}
FLAGS="-std=c++11 -fopenmp -O2 -fPIC -w"
for f in cpu_2dp cpu_2d cpu_3d Logging bin_io sys_file; do
  $CXX $FLAGS -I"$REF_SRC" -c "$REF_SRC/$f.cpp" -o "$OUT/$f.o"
done
$CXX $FLAGS -I"$REF_SRC" -c "$HERE/ref_harness.cpp" -o "$OUT/ref_harness.o"
$CXX -shared -fopenmp -o "$OUT/libref.so" "$OUT"/cpu_2dp.o "$OUT"/cpu_2d.o "$OUT"/cpu_3d.o \
  "$OUT"/Logging.o "$OUT"/bin_io.o "$OUT"/sys_file.o "$OUT"/ref_harness.o -lm
echo "build_ref: built $OUT/libref.so"
