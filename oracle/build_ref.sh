#!/bin/sh
# Compile the reference's own self-contained symbolic functions into
# oracle/_ref/libgfref.so, and its NASTRAN reader into
# oracle/_ref/libgfref_bdf.so, so the oracle and the product can be checked
# against them.
#
#   include/linalg/reordering_utils.h   compiled as is (TacsComputeRCMOrder)
#   include/linalg/bsr_utils.h          only its free functions, taken by line
#                                       range: the file as a whole includes the
#                                       sparse-utils headers, which are absent
#     :16       typedef index_t
#     :409-417  node_in_elem_conn
#     :430-476  get_row_col_ptrs
#     :655-670  getBandWidth
#     :678-807  computeILUk
#     :809-870  get_elem_ind_map
#
# The extracted text is written to a temporary directory and deleted after the
# compile; only the shared library lands in the output directory.
# usage: build_ref.sh <reference root> <output dir>
set -e
REF=${1:-/root/reference}
OUT=${2:-$(dirname "$0")/_ref}
B="$REF/include/linalg/bsr_utils.h"
if [ ! -f "$B" ] || [ ! -f "$REF/include/linalg/reordering_utils.h" ]; then
  echo "build_ref: reference tree not present at $REF, skipping"
  exit 0
fi
check() {  # line, expected text at the start of that line
  got=$(sed -n "${1}p" "$B")
  case "$got" in
    "$2"*) ;;
    *) echo "build_ref: $B:$1 is '$got', expected '$2'"; exit 1 ;;
  esac
}
check 16 "typedef int index_t;"
check 409 "__HOST_DEVICE__ bool node_in_elem_conn"
check 430 "__HOST__ void get_row_col_ptrs("
check 655 "__HOST__ int getBandWidth"
check 678 "__HOST__ void computeILUk"
check 809 "__HOST__ void get_elem_ind_map("
check 870 "}"
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
sed -n '16p;409,417p;430,476p;655,670p;678,807p;809,870p' "$B" > "$TMP/bsr_free_functions.inc"
cat > "$TMP/gfref.cpp" <<EOC
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <vector>
#include "$REF/include/cuda_utils.h"
#include "$REF/include/linalg/reordering_utils.h"
#include "bsr_free_functions.inc"
extern "C" {
int gfref_rcm(int n, const int* rowp, const int* cols, int root, int iters,
              int* rcm_vars) {
  return TacsComputeRCMOrder(n, rowp, cols, rcm_vars, root, iters);
}
int gfref_iluk(int n, int nnzb, const int* rowp, const int* cols, int lev,
               double fill, int** orowp, int** ocols, int** olevs) {
  int* rp = new int[n + 1];
  int* cp = new int[rowp[n] > 0 ? rowp[n] : 1];
  std::memcpy(rp, rowp, sizeof(int) * (n + 1));
  std::memcpy(cp, cols, sizeof(int) * rowp[n]);
  int* rp0 = rp;
  int* cp0 = cp;
  int* levs = nullptr;
  computeILUk(n, nnzb, rp, cp, lev, fill, &levs);
  delete[] rp0;
  delete[] cp0;
  *orowp = rp;
  *ocols = cp;
  *olevs = levs;
  return rp[n];
}
void gfref_elem_ind_map(int nelems, int nnodes, const int* conn, int npe,
                        int nnzb, int* rowp, int* cols, int* perm, int* out) {
  index_t* m = nullptr;
  get_elem_ind_map(nelems, nnodes, conn, npe, nnzb, rowp, cols, perm, m);
  std::memcpy(out, m, sizeof(int) * nelems * npe * npe);
  delete[] m;
}
int gfref_bandwidth(int n, int nnzb, int* rowp, int* cols) {
  return getBandWidth(n, nnzb, rowp, cols);
}
int gfref_row_col_ptrs(int nelems, int nnodes, const int* conn, int npe,
                       int** rowp, int** cols) {
  int nnzb = 0;
  index_t *rp = nullptr, *cp = nullptr;
  get_row_col_ptrs(nelems, nnodes, conn, npe, nnzb, rp, cp);
  *rowp = rp;
  *cols = cp;
  return nnzb;
}
void gfref_free(int* p) { delete[] p; }
}
EOC
mkdir -p "$OUT"
${CXX:-g++} -O2 -std=c++17 -fPIC -shared -w -I"$TMP" -o "$OUT/libgfref.so" "$TMP/gfref.cpp"
echo "build_ref: built $OUT/libgfref.so"

# the NASTRAN reader: TACSMeshLoader.h is self-contained (C library headers
# only) and is compiled where it lies, through the wrapper ref_bdf.cpp
L="$REF/include/mesh/TACSMeshLoader.h"
if [ -f "$L" ]; then
  ${CXX:-g++} -O2 -std=c++17 -fPIC -shared -w -I"$REF/include" \
    -o "$OUT/libgfref_bdf.so" "$(dirname "$0")/ref_bdf.cpp"
  echo "build_ref: built $OUT/libgfref_bdf.so"
fi
