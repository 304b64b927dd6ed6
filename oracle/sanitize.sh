#!/bin/bash
# AddressSanitizer + UndefinedBehaviorSanitizer run of the host code (test
# infrastructure; CPU only): builds the sanitized oracle library and the
# library's problem compiler (make sanitize), compiles a set of problem files
# with the sanitized compiler, and runs the CPU test suite against the
# sanitized oracle.  Exits non-zero on any sanitizer report or test failure.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(dirname "$HERE")
make -s -C "$HERE" sanitize
# libstdc++ preloaded too: Python does not link it, and ASan's exception
# interceptors need it present when ASan starts
ASAN_LIB="$(g++ -print-file-name=libasan.so) $(g++ -print-file-name=libstdc++.so.6)"
export ASAN_OPTIONS=detect_leaks=0:halt_on_error=1:abort_on_error=0
export UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
# problem files: 2-D and 3-D, every law, plasticity, curved surfaces,
# heterogeneous material and stress, fault input files, outputs and fronts
cd "$ROOT"
python3 - "$TMP" <<'PY'
import os, sys
sys.path.insert(0, "tests")
sys.path.insert(0, ".")
from cases import STZ_PARAMS, curved_surfaces, reference_example, write_case
from test_oracle_law_pins import file_cases
d = sys.argv[1]
paths = [reference_example(d, "test2d", 5, d + "/data/"), reference_example(d, "tpv4", 3, d + "/data/"),
         write_case(d, "m3", mode=3, nx=41, ny=21, dx=0.1),
         write_case(d, "st", nx=41, ny=21, dx=0.1, law="stz", stz=dict(STZ_PARAMS, chi0=0.07)),
         write_case(d, "pl", nx=41, ny=21, dx=0.1, material="plastic",
                    plastic=dict(mu=0.68, c=0., beta=0.35, eta=1.6), s0=(-100., 52., 0., -120., 0., -55.)),
         write_case(d, "cv", nx=41, ny=21, dx=0.1, surf=curved_surfaces(d, 41, 21, 0.1, amp=0.1)),
         write_case(d, "s3", ndim=3, nx=21, ny=13, nz=15, dx=0.1, law="stz", stz=dict(STZ_PARAMS, chi0=0.07))]
for nd in (2, 3):
    for w in ("load", "sw", "stz"):
        paths += list(file_cases(d, nd, w))
open(os.path.join(d, "list.txt"), "w").write("\n".join(paths))
PY
echo "== problem compiler (setup.cpp) on $(wc -l < "$TMP/list.txt") problem files"
xargs "$HERE/_build/asan/setup_check" < "$TMP/list.txt" > "$TMP/setup.log" 2>&1 || { cat "$TMP/setup.log"; exit 1; }
tail -3 "$TMP/setup.log"
echo "== CPU test suite against the sanitized oracle"
LD_PRELOAD="$ASAN_LIB" FDF_ORACLE_LIB="$HERE/_build/asan/liboracle.so" OMP_NUM_THREADS=4 \
  python3 -m pytest "$ROOT/tests" -q -m "not gpu" -p no:cacheprovider -x
echo "== sanitizers: no reports"
