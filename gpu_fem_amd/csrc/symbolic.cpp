// Host symbolic phase of the product library: node graph, reorderings,
// pattern permutation, ILU(k) fill, element->block map, transpose map and
// the level schedules used by the device ILU and triangular sweeps.
//
// These are the product's own implementations (array-based; O(nnz) work,
// and for ILU(k) one step per (row, pivot, U entry of the pivot row)); they
// must reproduce the reference's patterns bit for bit:
//   gf_node_graph       bsr_utils.h:419-476
//   gf_rcm_order        reordering_utils.h:189-291
//   gf_pattern_permute  gmres.cu:242-252 (BSRMatApplyPerm)
//   gf_iluk_symbolic    bsr_utils.h:678-807 (computeILUk)
//   gf_elem_ind_map     bsr_utils.h:809-870
//   gf_transpose_map    bsr_utils.h:157-211
//   gf_bandwidth        bsr_utils.h:655-670
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

#include "symbolic.h"

namespace {
int* host_copy(const std::vector<int>& v) {
  int* p = (int*)std::malloc(sizeof(int) * (v.empty() ? 1 : v.size()));
  if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(int) * v.size());
  return p;
}
}  // namespace

namespace gf {

int node_graph(int nelems, int nnodes, int npe, const int* conn,
               std::vector<int>& rowp, std::vector<int>& cols) {
  // node -> element incidence (CSR), then per node the sorted union of the
  // incident elements' nodes.
  std::vector<int> eptr(nnodes + 1, 0), elist((size_t)nelems * npe);
  for (size_t i = 0; i < (size_t)nelems * npe; i++) {
    if (conn[i] < 0 || conn[i] >= nnodes) return GF_ERR_INVALID;
    eptr[conn[i] + 1]++;
  }
  for (int i = 0; i < nnodes; i++) eptr[i + 1] += eptr[i];
  std::vector<int> fillp(eptr.begin(), eptr.end() - 1);
  for (int e = 0; e < nelems; e++)
    for (int a = 0; a < npe; a++) elist[fillp[conn[npe * e + a]]++] = e;
  rowp.assign(nnodes + 1, 0);
  cols.clear();
  cols.reserve((size_t)nnodes * 10);
  std::vector<int> buf;
  for (int i = 0; i < nnodes; i++) {
    buf.clear();
    for (int k = eptr[i]; k < eptr[i + 1]; k++) {
      const int* ec = &conn[(size_t)npe * elist[k]];
      buf.insert(buf.end(), ec, ec + npe);
    }
    if (buf.empty()) buf.push_back(i);  // isolated node keeps its diagonal
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
    cols.insert(cols.end(), buf.begin(), buf.end());
    rowp[i + 1] = (int)cols.size();
  }
  return GF_OK;
}

int rcm_order(int n, const int* rowp, const int* cols, int root, int iters,
              int* rcm_vars) {
  if (n <= 0 || root < 0 || root >= n) return GF_ERR_INVALID;
  if (iters < 1) iters = 1;
  std::vector<int> order(n);
  for (int it = 0; it < iters; it++) {
    std::fill(rcm_vars, rcm_vars + n, -1);
    int head = 0, tail = 0, numbered = 0, scan = 0;
    while (numbered < n) {
      if (head == tail) {
        // new component: the given root first, then the lowest unnumbered
        int r = root;
        if (rcm_vars[r] >= 0) {
          while (scan < n && rcm_vars[scan] >= 0) scan++;
          r = scan;
        }
        rcm_vars[r] = numbered++;
        order[tail++] = r;
      }
      // one breadth-first sweep over the current queue segment
      int end = tail;
      for (; head < end; head++) {
        int node = order[head];
        for (int j = rowp[node]; j < rowp[node + 1]; j++) {
          int nb = cols[j];
          if (rcm_vars[nb] < 0) {
            rcm_vars[nb] = numbered++;
            order[tail++] = nb;
          }
        }
      }
    }
    for (int i = 0; i < n; i++) rcm_vars[order[i]] = n - 1 - i;
    root = rcm_vars[0];
  }
  return GF_OK;
}

int pattern_permute(int n, const int* rowp, const int* cols, const int* perm,
                    std::vector<int>& prowp, std::vector<int>& pcols) {
  std::vector<int> iperm(n, -1);
  for (int i = 0; i < n; i++) {
    if (perm[i] < 0 || perm[i] >= n || iperm[perm[i]] >= 0)
      return GF_ERR_INVALID;
    iperm[perm[i]] = i;
  }
  prowp.assign(n + 1, 0);
  for (int r = 0; r < n; r++) {
    int o = iperm[r];
    prowp[r + 1] = prowp[r] + (rowp[o + 1] - rowp[o]);
  }
  pcols.resize(prowp[n]);
  for (int r = 0; r < n; r++) {
    int o = iperm[r];
    int* out = &pcols[prowp[r]];
    int len = rowp[o + 1] - rowp[o];
    for (int k = 0; k < len; k++) out[k] = perm[cols[rowp[o] + k]];
    std::sort(out, out + len);
  }
  return GF_OK;
}

namespace {
struct int2s {
  int row, lev;
};
}  // namespace

int iluk_symbolic(int n, const int* rowp, const int* cols, int lev_fill,
                  std::vector<int>& frowp, std::vector<int>& fcols,
                  std::vector<int>* flevs) {
  // Row-by-row level-of-fill factorisation.  The active row's levels live in
  // a dense array (a row stamp per column marks it present in row i), its
  // pivots (columns left of the diagonal, existing or created by fill) come
  // off a min-heap in ascending order, and the columns right of the
  // diagonal are collected unsorted and sorted once when the row is
  // emitted.  A pivot k merges the U part of row k with
  //   lev(i, c) = min(lev(i, c), lev(i, k) + lev(k, c) + 1),
  // keeping a new column only if that level is at most lev_fill; every
  // pivot below c is applied before c itself is taken as a pivot, so its
  // level is final by then.  The work is one constant-time update per
  // (row, pivot, U entry of the pivot row), with no sorted-list walk.
  frowp.assign(n + 1, 0);
  fcols.clear();
  std::vector<int> levs, udiag(n);
  fcols.reserve((size_t)rowp[n] * 2);
  levs.reserve((size_t)rowp[n] * 2);
  // per column: {row stamp, level}, side by side so an update touches one
  // cache line
  std::vector<int2s> st(n, int2s{-1, 0});
  std::vector<int> heap, ucols;
  auto heap_push = [&](int c) {
    heap.push_back(c);
    std::push_heap(heap.begin(), heap.end(), std::greater<int>());
  };
  for (int i = 0; i < n; i++) {
    heap.clear();
    ucols.clear();
    st[i] = int2s{i, 0};
    for (int k = rowp[i]; k < rowp[i + 1]; k++) {
      const int c = cols[k];
      if (c == i) continue;
      st[c] = int2s{i, 0};
      if (c < i)
        heap.push_back(c);
      else
        ucols.push_back(c);
    }
    std::make_heap(heap.begin(), heap.end(), std::greater<int>());
    while (!heap.empty()) {
      std::pop_heap(heap.begin(), heap.end(), std::greater<int>());
      const int k = heap.back();
      heap.pop_back();
      const int lk = st[k].lev;
      fcols.push_back(k);
      levs.push_back(lk);
      const int* uc = fcols.data() + udiag[k] + 1;
      const int* ul = levs.data() + udiag[k] + 1;
      const int nu = frowp[k + 1] - (udiag[k] + 1);
      for (int t = 0; t < nu; t++) {
        const int c = uc[t];
        const int nl = lk + ul[t] + 1;
        int2s& e = st[c];
        if (e.row == i) {
          if (e.lev > nl) e.lev = nl;
        } else if (nl <= lev_fill) {
          e = int2s{i, nl};
          if (c < i)
            heap_push(c);
          else
            ucols.push_back(c);
        }
      }
    }
    udiag[i] = (int)fcols.size();
    fcols.push_back(i);
    levs.push_back(0);
    std::sort(ucols.begin(), ucols.end());
    for (int c : ucols) {
      fcols.push_back(c);
      levs.push_back(st[c].lev);
    }
    frowp[i + 1] = (int)fcols.size();
  }
  if (flevs) flevs->swap(levs);
  return GF_OK;
}

int elem_ind_map(int nelems, int npe, const int* conn, const int* rowp,
                 const int* cols, const int* perm, int* map) {
  int npe2 = npe * npe;
  for (int e = 0; e < nelems; e++) {
    const int* lc = &conn[(size_t)npe * e];
    for (int a = 0; a < npe; a++) {
      int row = perm[lc[a]];
      const int* b0 = cols + rowp[row];
      const int* b1 = cols + rowp[row + 1];
      for (int b = 0; b < npe; b++) {
        int col = perm[lc[b]];
        const int* it = std::lower_bound(b0, b1, col);
        if (it == b1 || *it != col) return GF_ERR_INVALID;
        map[(size_t)npe2 * e + npe * a + b] = (int)(it - cols);
      }
    }
  }
  return GF_OK;
}

int transpose_map(int n, const int* rowp, const int* cols,
                  std::vector<int>& trowp, std::vector<int>& tcols,
                  std::vector<int>& tmap) {
  int nnz = rowp[n];
  trowp.assign(n + 1, 0);
  for (int k = 0; k < nnz; k++) trowp[cols[k] + 1]++;
  for (int i = 0; i < n; i++) trowp[i + 1] += trowp[i];
  tcols.resize(nnz);
  tmap.resize(nnz);
  std::vector<int> pos(trowp.begin(), trowp.end() - 1);
  for (int r = 0; r < n; r++)
    for (int k = rowp[r]; k < rowp[r + 1]; k++) {
      int t = pos[cols[k]]++;
      tcols[t] = r;
      tmap[t] = k;
    }
  return GF_OK;
}

int bandwidth(int n, const int* rowp, const int* cols) {
  int bw = 0;
  for (int i = 0; i < n; i++)
    for (int k = rowp[i]; k < rowp[i + 1]; k++)
      bw = std::max(bw, std::abs(i - cols[k]));
  return bw;
}

void level_schedule(int n, const int* rowp, const int* cols, bool lower,
                    std::vector<int>& level_ptr, std::vector<int>& level_rows) {
  std::vector<int> lev(n, 0);
  int nlev = 0;
  if (lower) {
    for (int i = 0; i < n; i++) {
      int l = 0;
      for (int k = rowp[i]; k < rowp[i + 1] && cols[k] < i; k++)
        l = std::max(l, lev[cols[k]] + 1);
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
  } else {
    for (int i = n - 1; i >= 0; i--) {
      int l = 0;
      for (int k = rowp[i + 1] - 1; k >= rowp[i] && cols[k] > i; k--)
        l = std::max(l, lev[cols[k]] + 1);
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
  }
  level_ptr.assign(nlev + 1, 0);
  for (int i = 0; i < n; i++) level_ptr[lev[i] + 1]++;
  for (int l = 0; l < nlev; l++) level_ptr[l + 1] += level_ptr[l];
  level_rows.resize(n);
  std::vector<int> pos(level_ptr.begin(), level_ptr.end() - 1);
  // rows inside a level keep ascending (lower) / descending (upper) order
  if (lower) {
    for (int i = 0; i < n; i++) level_rows[pos[lev[i]]++] = i;
  } else {
    for (int i = n - 1; i >= 0; i--) level_rows[pos[lev[i]]++] = i;
  }
}

void element_coloring(int nelems, int nnodes, int npe, const int* conn,
                      std::vector<int>& color_ptr,
                      std::vector<int>& color_elems) {
  // node -> elements
  std::vector<int> nptr(nnodes + 1, 0), nel((size_t)nelems * npe);
  for (size_t k = 0; k < (size_t)nelems * npe; k++) nptr[conn[k] + 1]++;
  for (int i = 0; i < nnodes; i++) nptr[i + 1] += nptr[i];
  {
    std::vector<int> pos(nptr.begin(), nptr.end() - 1);
    for (int e = 0; e < nelems; e++)
      for (int a = 0; a < npe; a++) nel[pos[conn[(size_t)npe * e + a]]++] = e;
  }
  std::vector<int> color(nelems, -1), mark;
  int ncolors = 0;
  for (int e = 0; e < nelems; e++) {
    mark.assign(ncolors + 1, 0);
    for (int a = 0; a < npe; a++) {
      const int nd = conn[(size_t)npe * e + a];
      for (int k = nptr[nd]; k < nptr[nd + 1]; k++) {
        const int c = color[nel[k]];
        if (c >= 0) mark[c] = 1;
      }
    }
    int c = 0;
    while (mark[c]) c++;
    color[e] = c;
    ncolors = std::max(ncolors, c + 1);
  }
  color_ptr.assign(ncolors + 1, 0);
  for (int e = 0; e < nelems; e++) color_ptr[color[e] + 1]++;
  for (int c = 0; c < ncolors; c++) color_ptr[c + 1] += color_ptr[c];
  color_elems.resize(nelems);
  std::vector<int> pos(color_ptr.begin(), color_ptr.end() - 1);
  for (int e = 0; e < nelems; e++) color_elems[pos[color[e]]++] = e;
}

}  // namespace gf

// ------------------------------------------------------------------ C ABI
extern "C" {

void gf_free(void* p) { std::free(p); }

int gf_node_graph(int nelems, int nnodes, int npe, const int* conn, int** rowp,
                  int** cols, int* nnzb) {
  if (nelems < 0 || nnodes <= 0 || npe <= 0 || !conn || !rowp || !cols)
    return GF_ERR_INVALID;
  std::vector<int> rp, cl;
  GF_TRY(gf::node_graph(nelems, nnodes, npe, conn, rp, cl));
  *rowp = host_copy(rp);
  *cols = host_copy(cl);
  if (nnzb) *nnzb = rp[nnodes];
  return (*rowp && *cols) ? GF_OK : GF_ERR_NOMEM;
}

int gf_rcm_order(int n, const int* rowp, const int* cols, int root, int iters,
                 int* rcm_vars) {
  if (!rowp || !cols || !rcm_vars) return GF_ERR_INVALID;
  return gf::rcm_order(n, rowp, cols, root, iters, rcm_vars);
}

int gf_pattern_permute(int n, const int* rowp, const int* cols,
                       const int* perm, int** prowp, int** pcols) {
  if (n <= 0 || !rowp || !cols || !perm || !prowp || !pcols)
    return GF_ERR_INVALID;
  std::vector<int> rp, cl;
  GF_TRY(gf::pattern_permute(n, rowp, cols, perm, rp, cl));
  *prowp = host_copy(rp);
  *pcols = host_copy(cl);
  return (*prowp && *pcols) ? GF_OK : GF_ERR_NOMEM;
}

int gf_iluk_symbolic(int n, const int* rowp, const int* cols, int lev_fill,
                     int** frowp, int** fcols, int* nnzb) {
  if (n <= 0 || lev_fill < 0 || !rowp || !cols || !frowp || !fcols)
    return GF_ERR_INVALID;
  std::vector<int> rp, cl;
  GF_TRY(gf::iluk_symbolic(n, rowp, cols, lev_fill, rp, cl, nullptr));
  *frowp = host_copy(rp);
  *fcols = host_copy(cl);
  if (nnzb) *nnzb = rp[n];
  return (*frowp && *fcols) ? GF_OK : GF_ERR_NOMEM;
}

int gf_elem_ind_map(int nelems, int npe, const int* conn, const int* rowp,
                    const int* cols, const int* perm, int* map) {
  if (nelems < 0 || npe <= 0 || !conn || !rowp || !cols || !perm || !map)
    return GF_ERR_INVALID;
  return gf::elem_ind_map(nelems, npe, conn, rowp, cols, perm, map);
}

int gf_transpose_map(int n, const int* rowp, const int* cols, int** trowp,
                     int** tcols, int** tmap) {
  if (n <= 0 || !rowp || !cols || !trowp || !tcols || !tmap)
    return GF_ERR_INVALID;
  std::vector<int> tr, tc, tm;
  GF_TRY(gf::transpose_map(n, rowp, cols, tr, tc, tm));
  *trowp = host_copy(tr);
  *tcols = host_copy(tc);
  *tmap = host_copy(tm);
  return GF_OK;
}

int gf_bandwidth(int n, const int* rowp, const int* cols) {
  if (n <= 0 || !rowp || !cols) return GF_ERR_INVALID;
  return gf::bandwidth(n, rowp, cols);
}

}  // extern "C"
