#!/bin/bash
# Builds the parts of the reference that compile from their own sources, into oracle/_ref/
# (git-ignored). Only the formula evaluator does: src/eqEdit.cpp + include/eqEdit.h. The
# solver itself needs the gmsh SDK, Eigen, VTK and FFTW, which are not available.
set -e
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
mkdir -p "$HERE/_ref"
g++ -O1 -std=c++17 -w -I"$REF/include" "$REF/src/eqEdit.cpp" "$HERE/eqedit_ref_driver.cpp" -o "$HERE/_ref/eqedit_ref"
echo "built $HERE/_ref/eqedit_ref"
