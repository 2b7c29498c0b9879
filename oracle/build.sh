#!/bin/sh
# Builds the CPU oracle (test infrastructure) twice: double and float.
# Usage: sh oracle/build.sh   (outputs oracle/liboracle_f64.so, oracle/liboracle_f32.so)
set -e
here=$(cd "$(dirname "$0")" && pwd)
for p in f64:double f32:float; do
  tag=${p%%:*}; real=${p#*:}
  gcc -O2 -std=gnu11 -fopenmp -fPIC -shared -Wall -Wextra -Wno-unused-parameter \
      -DREAL=$real -o "$here/liboracle_$tag.so" "$here/oracle.c" -lm
done
