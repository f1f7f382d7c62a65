#!/bin/sh
# Builds oracle/_ref/ref_vectors: the reference's own host-only C++ (row-mix
# recipe, RNG draws, instance loader, sector layouts and the exact CPU apply
# of the eliminated operator) compiled straight from the reference sources,
# plus this repository's driver oracle/ref_vectors.cpp, which calls those
# functions and prints golden vectors.  Test infrastructure only.
#
# The reference's translation unit cuda/cuprym_cyclic.cu also holds CUDA
# kernels and device code, so only its host-only line ranges are fed to g++
# (on stdin, nothing is copied into this repository), together with the
# plain-C++ structs of cuda/cyclic_cuda_helpers.h and the header
# cuda/matrices.h (pure C++).  Every range is standard C++17; nothing is
# stubbed or replaced.
#   cuda/cyclic_cuda_helpers.h:15-74   descriptor/incidence/row-mix structs, CyclicPhiHostData
#   cuda/cuprym_cyclic.cu:29-885       typedefs, instance structs, subsets, JSON loader,
#                                      elimination checks, character maps, incidences,
#                                      sector/tensor layouts, operator host data
#   cuda/cuprym_cyclic.cu:1009-1421    accumulation guards, RNG, row-mix plan and hash,
#                                      sidecar, deterministic blocks
#   cuda/cuprym_cyclic.cu:1433-1637    dense helpers, CPU row mixing, CPU forward differential
#   cuda/cuprym_cyclic.cu:1685-1799    CPU target preconditioner, elimination, eliminated forward
# The CPU transpose differential (:1638-1683) and everything after it call
# cyclic_add_mod, a __host__ __device__ function of cyclic_cuda_helpers.h that
# needs the CUDA headers, so the reference's backward half is not built: its
# forward F (Dc x) and P^T Dr P are, and F^T is pinned by transpose identities.
set -e
REF=${PGW_REFERENCE:-/root/reference}
here=$(cd "$(dirname "$0")" && pwd)
out="$here/_ref"
if [ ! -f "$REF/cuda/cuprym_cyclic.cu" ]; then
    echo "ref_build: no reference sources at $REF; nothing built" >&2
    exit 0
fi
mkdir -p "$out"
{
    sed -n '15,74p' "$REF/cuda/cyclic_cuda_helpers.h"
    sed -n '29,885p' "$REF/cuda/cuprym_cyclic.cu"
    sed -n '1009,1421p' "$REF/cuda/cuprym_cyclic.cu"
    sed -n '1433,1637p' "$REF/cuda/cuprym_cyclic.cu"
    sed -n '1685,1799p' "$REF/cuda/cuprym_cyclic.cu"
    cat "$here/ref_vectors.cpp"
} | g++ -O2 -std=c++17 -w -x c++ \
    -include "$REF/cuda/matrices.h" -include unordered_map -include cctype \
    -o "$out/ref_vectors.tmp" -
mv "$out/ref_vectors.tmp" "$out/ref_vectors"
echo "ref_build: built $out/ref_vectors"
