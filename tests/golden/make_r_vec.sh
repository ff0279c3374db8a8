#!/bin/sh
# Regenerates r_vec_flang.txt with the image's Fortran compiler (amdflang).
set -e
cd "$(dirname "$0")"
tmp=$(mktemp -d)
/opt/rocm/bin/amdflang -O2 r_vec_flang.f90 -o "$tmp/r_vec_flang"
"$tmp/r_vec_flang" > r_vec_flang.txt
rm -rf "$tmp"
