#!/bin/bash
# Attempts to compile the reference's hot-path headers into oracle/_ref/ as a
# second checker.  Every reference header includes
# include/cdet/utility/types.h, which needs boost/multiprecision (not present
# in this image and not substituted here), so the reference is unbuildable and
# this recipe stops after the probe.  Output (if any) goes to oracle/_ref/.
set -u
HERE="$(cd "$(dirname "$0")" && pwd)"
REF=/root/reference
mkdir -p "$HERE/_ref"
if [ ! -d "$REF" ]; then
  echo "reference not present; nothing to build"
  exit 0
fi
if echo '#include <boost/multiprecision/cpp_dec_float.hpp>' | g++ -x c++ -fsyntax-only - 2>/dev/null; then
  g++ -O3 -std=c++20 -fPIC -shared -I"$REF/include" -o "$HERE/_ref/libcdet_ref_probe.so" \
      -x c++ - <<'CPP'
#include "cdet/fast_principle_minors.h"
#include "cdet/integrand_replica_randomly_sampled.h"
extern "C" int cdet_ref_probe() { return 1; }
CPP
  echo "built $HERE/_ref/libcdet_ref_probe.so"
else
  echo "reference unbuildable here: boost/multiprecision/cpp_dec_float.hpp is missing (include/cdet/utility/types.h:4)" \
    > "$HERE/_ref/UNBUILDABLE.txt"
  cat "$HERE/_ref/UNBUILDABLE.txt"
fi
exit 0
