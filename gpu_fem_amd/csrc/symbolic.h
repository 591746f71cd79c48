// Host symbolic phase (see symbolic.cpp for the reference functions each one
// reproduces).
#pragma once
#include <vector>

#include "../../include/gpufem_amd.h"

#define GF_TRY_H(expr)            \
  do {                            \
    int _st = (expr);             \
    if (_st != GF_OK) return _st; \
  } while (0)
#ifndef GF_TRY
#define GF_TRY GF_TRY_H
#endif

namespace gf {
int node_graph(int nelems, int nnodes, int npe, const int* conn,
               std::vector<int>& rowp, std::vector<int>& cols);
int rcm_order(int n, const int* rowp, const int* cols, int root, int iters,
              int* rcm_vars);
int amd_order(int n, const int* rowp, const int* cols, int* perm);
int pattern_permute(int n, const int* rowp, const int* cols, const int* perm,
                    std::vector<int>& prowp, std::vector<int>& pcols);
int iluk_symbolic(int n, const int* rowp, const int* cols, int lev_fill,
                  std::vector<int>& frowp, std::vector<int>& fcols,
                  std::vector<int>* flevs);
int elem_ind_map(int nelems, int npe, const int* conn, const int* rowp,
                 const int* cols, const int* perm, int* map);
int transpose_map(int n, const int* rowp, const int* cols,
                  std::vector<int>& trowp, std::vector<int>& tcols,
                  std::vector<int>& tmap);
int bandwidth(int n, const int* rowp, const int* cols);
// rows grouped by dependency level for the lower (L) or upper (U) sweep
void level_schedule(int n, const int* rowp, const int* cols, bool lower,
                    std::vector<int>& level_ptr, std::vector<int>& level_rows);
// Greedy element colouring: no two elements of a colour share a node.
// Elements are visited in index order and take the smallest colour unused by
// the elements already coloured around their nodes.  Returns the colour
// offsets (color_ptr) and the elements in colour-major, index order
// (color_elems).
void element_coloring(int nelems, int nnodes, int nodes_per_elem,
                      const int* conn, std::vector<int>& color_ptr,
                      std::vector<int>& color_elems);
}  // namespace gf
