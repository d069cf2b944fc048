#!/usr/bin/env bash
# Builds the reference DetNQS C++ core (hot-path translation units only, read in
# place from /root/reference/csrc) plus oracle/ref_capi.cpp into
# oracle/_ref/libdetnqs_ref.so.  Used only as a test checker and as the CPU
# baseline; never shipped in the product and never committed (oracle/_ref/ is
# git-ignored).  det_enum.cpp is left out: it needs <format>, absent from
# this image's g++ 11.4.
#
# Flags mirror the reference's Release build (csrc/CMakeLists.txt:13-37:
# C++20, OpenMP, PIC, no -march), i.e. no FMA contraction on x86-64.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
ref="${DETNQS_REFERENCE:-/root/reference}/csrc"
out="$here/_ref"
if [ ! -d "$ref" ]; then
  echo "build_ref: reference sources not present at $ref; skipping" >&2
  exit 0
fi
mkdir -p "$out"
srcs=(
  "$ref/src/determinant/det_ops.cpp"
  "$ref/src/determinant/det_batch.cpp"
  "$ref/src/determinant/det_space.cpp"
  "$ref/src/integral/integral_mo.cpp"
  "$ref/src/integral/hb_table.cpp"
  "$ref/src/hamiltonian/ham_eval.cpp"
  "$ref/src/hamiltonian/ham_utils.cpp"
  "$ref/src/hamiltonian/build_ham.cpp"
  "$ref/src/hamiltonian/local_conn.cpp"
  "$ref/src/hamiltonian/ham_eff.cpp"
)
g++ -std=c++20 -O3 -fopenmp -fPIC -shared -DNDEBUG -ffp-contract=off \
  -I"$ref" "$here/ref_capi.cpp" "${srcs[@]}" -o "$out/libdetnqs_ref.so.tmp"
mv "$out/libdetnqs_ref.so.tmp" "$out/libdetnqs_ref.so"
echo "build_ref: wrote $out/libdetnqs_ref.so"
